"""GPU: the wave-cooperative self-query search (float trees, every point
queried) against the oracle and against the per-lane search.

Queries given as an index list and normals for explicit rows run the per-lane
search (one traversal per query); queries=None and normals of every row run
the cooperative one. Both must give the same rows and the same normals bit
for bit: the neighbour set and its order are defined by the (distance,
original index) key, whichever search finds it."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mixed_density_cloud(seed):
    """Dense clusters, sparse shells, isolated points and exact duplicates:
    waves whose 64 Morton-consecutive points have very different k-th
    distances, and waves that straddle far-apart clusters."""
    rng = np.random.default_rng(seed)
    parts = []
    for c in range(6):
        centre = rng.uniform(-50, 50, 3)
        parts.append(centre + rng.normal(scale=10 ** rng.uniform(-3, 0), size=(rng.integers(2000, 8000), 3)))
    parts.append(rng.uniform(-200, 200, (300, 3)))                   # isolated points
    shell = rng.normal(size=(3000, 3))
    parts.append(80 * shell / np.linalg.norm(shell, axis=1, keepdims=True))  # sparse shell
    pts = np.concatenate(parts).astype(np.float32)
    dup = pts[rng.choice(len(pts), 500)]
    pts = np.concatenate([pts, dup, dup[:50]])                       # exact duplicates
    return pts[rng.permutation(len(pts))]


@pytest.mark.parametrize("k", [1, 4, 7, 16, 24, 32])
def test_coop_knn_mixed_density_vs_oracle(gpu, k):
    from pptk_amd import capi
    pts = _mixed_density_cloud(7)
    t = capi.Tree.build(pts)
    gi, gc = capi.knn(t, None, k)
    wi, wc = O.OracleTree(pts).knn(None, k)
    assert np.array_equal(gc, wc)
    assert np.array_equal(gi, wi)


@pytest.mark.parametrize("dmax", [0.05, 0.5, 5.0])
def test_coop_knn_dmax_vs_oracle(gpu, dmax):
    from pptk_amd import capi
    pts = _mixed_density_cloud(8)
    t = capi.Tree.build(pts)
    d32 = float(np.float32(dmax))
    gi, gc = capi.knn(t, None, 12, dmax=d32)
    wi, wc = O.OracleTree(pts).knn(None, 12, dmax=d32)
    assert np.array_equal(gc, wc)
    assert np.array_equal(gi, wi)


@pytest.mark.parametrize("leaf", [1, 5, 32, 256])
def test_coop_knn_equals_per_lane(gpu, leaf):
    """queries=None (cooperative) against the same queries as an index list
    (per-lane), on a lidar-like scene, with distances."""
    from pptk_amd import capi, scenes
    pts = scenes.lidar_scene(200_000, seed=5)
    t = capi.Tree.build(pts, leaf=leaf)
    a = capi.knn(t, None, 16, with_dist=True)
    b = capi.knn(t, np.arange(len(pts)), 16, with_dist=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("k", [3, 16, 30])
def test_coop_normals_equal_per_lane(gpu, k):
    """Normals of every row (cooperative search) against normals of the same
    rows listed explicitly (per-lane search): identical bits."""
    from pptk_amd import capi
    pts = _mixed_density_cloud(9)
    t = capi.Tree.build(pts)
    rows = np.arange(len(pts), dtype=np.int32)
    a = capi.normals(t, k, np.inf, num_eigen=3)
    b = capi.normals(t, k, np.inf, rows=rows, num_eigen=3)
    assert np.array_equal(a[0], b[0], equal_nan=True)
    assert np.array_equal(a[1], b[1], equal_nan=True)
    assert np.array_equal(a[2], b[2])


def test_coop_normals_radius_equal_per_lane(gpu):
    from pptk_amd import capi
    pts = _mixed_density_cloud(10)
    t = capi.Tree.build(pts)
    rows = np.arange(len(pts), dtype=np.int32)
    r = float(np.float32(0.3))
    a = capi.normals(t, 16, r)
    b = capi.normals(t, 16, r, rows=rows)
    assert np.array_equal(a[0], b[0], equal_nan=True)
    assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("dim", [2, 4])
def test_coop_knn_other_dims(gpu, dim):
    from pptk_amd import capi
    rng = np.random.default_rng(dim)
    pts = np.concatenate([rng.random((20_000, dim)), 30 + rng.random((300, dim)) * 50]).astype(np.float32)
    t = capi.Tree.build(pts)
    gi, gc = capi.knn(t, None, 9)
    wi, wc = O.OracleTree(pts).knn(None, 9)
    assert np.array_equal(gc, wc)
    assert np.array_equal(gi, wi)


def test_coop_batch_mixed_clouds(gpu):
    """A batch whose waves span several clouds of very different scales
    (segments per cloud and per outlier class), against the oracle per cloud."""
    from pptk_amd import capi
    import torch
    rng = np.random.default_rng(12)
    clouds = []
    for i in range(60):
        n = int(rng.integers(1, 400)) if i % 3 else int(rng.integers(500, 3000))
        scale = 10 ** rng.uniform(-3, 3)
        clouds.append((rng.normal(size=(n, 3)) * scale).astype(np.float32))
    clouds[5] = np.zeros((0, 3), np.float32)
    sizes = np.array([len(c) for c in clouds], np.int64)
    allp = np.concatenate(clouds)
    N = len(allp)
    k = 10
    dev = torch.device("cuda:0")
    pts_d = torch.from_numpy(allp).to(dev)
    idx_d = torch.empty((max(N, 1), k), dtype=torch.int32, device=dev)
    cnt_d = torch.empty(max(N, 1), dtype=torch.int32, device=dev)
    plan = capi.Batch(sizes, k=k, normals_k=0, leaf=8, query_block=64)
    torch.cuda.synchronize()
    plan.run(pts_d.data_ptr(), idx_d.data_ptr(), cnt_d.data_ptr(), None, None)
    torch.cuda.synchronize()
    gi, gc = idx_d.cpu().numpy(), cnt_d.cpu().numpy()
    plan.free()
    off = 0
    for c in clouds:
        n = len(c)
        if n:
            wi, wc = O.OracleTree(c).knn(None, k)
            assert np.array_equal(gc[off:off + n], wc)
            assert np.array_equal(gi[off:off + n], wi)
        off += n
