"""GPU: batched clouds (pptk_batch_*) against the single-cloud path and the
oracle.

Each cloud of a batch must give exactly what the single-cloud calls give on
that cloud alone: kdtree._query(None, k) self-query rows (indices local to
the cloud) and estimate_normals(k, r) normals. Both are compared bit for bit
(same neighbour sets, same summation order), and a sample of clouds is also
checked against the oracle's kNN and FP64 normals. The hip-graph replay is
checked against an eager run, on two different inputs through one plan.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _edge_clouds(rng, dim=3, dtype=np.float32):
    """Clouds covering the edge cases: empty, 1 and 2 points, fewer than k+1
    points, exactly k+1, duplicates, NaN rows, an all-NaN cloud, a flat
    cloud, and log-uniform sizes up to 20k."""
    clouds = []
    for n in [0, 1, 2, 5, 16, 17, 33]:
        clouds.append(rng.uniform(0, 1, (n, dim)))
    dup = np.repeat(rng.uniform(0, 1, (40, dim)), 3, axis=0)
    clouds.append(dup)
    nan = rng.uniform(0, 1, (500, dim))
    nan[rng.choice(500, 37, replace=False), rng.integers(0, dim, 37)] = np.nan
    nan[3] = np.inf
    clouds.append(nan)
    clouds.append(np.full((20, dim), np.nan))
    flat = rng.uniform(0, 1, (3000, dim))
    flat[:, -1] = 0.25
    clouds.append(flat)
    same = np.zeros((70, dim))
    clouds.append(same)
    sizes = np.exp(rng.uniform(np.log(50), np.log(20_000), 40)).astype(np.int64)
    for s in sizes:
        c = rng.normal(0, 1, dim)
        clouds.append(c + rng.normal(0, rng.uniform(0.1, 3), (int(s), dim)))
    order = rng.permutation(len(clouds))
    return [np.ascontiguousarray(clouds[i], dtype=dtype) for i in order]


def run_batch(clouds, dim, dtype, k, nk, dmax=np.inf, leaf=0, use_graph=True, qblock=0,
              plan=None, want_eval=True):
    """Runs a batch over device copies; returns (idx, cnt, nrm, ev, plan)."""
    from pptk_amd import capi
    torch = _torch()
    sizes = [len(c) for c in clouds]
    if plan is None:
        plan = capi.Batch(sizes, dim=dim, dtype=dtype, k=k, normals_k=nk, dmax=dmax, leaf=leaf,
                          use_graph=use_graph, query_block=qblock)
    n = plan.num_points
    tdt = torch.float64 if np.dtype(dtype) == np.float64 else torch.float32
    allp = np.concatenate(clouds).reshape(-1, dim) if n else np.zeros((0, dim), dtype)
    if not hasattr(plan, "_bufs"):
        dev = torch.device("cuda:0")
        plan._bufs = {
            "pts": torch.empty((max(n, 1), dim), dtype=tdt, device=dev),
            "idx": torch.empty((max(n, 1), max(k, 1)), dtype=torch.int32, device=dev),
            "cnt": torch.empty(max(n, 1), dtype=torch.int32, device=dev),
            "nrm": torch.empty((max(n, 1), 3), dtype=tdt, device=dev),
            "ev": torch.empty(max(n, 1), dtype=tdt, device=dev),
        }
    b = plan._bufs
    if n:
        b["pts"][:n].copy_(torch.from_numpy(np.ascontiguousarray(allp, dtype=dtype)))
    b["idx"].fill_(-7)
    b["nrm"].fill_(7.0)
    torch.cuda.synchronize()
    plan.run(b["pts"].data_ptr(),
             b["idx"].data_ptr() if k else None,
             b["cnt"].data_ptr() if k else None,
             b["nrm"].data_ptr() if nk else None,
             b["ev"].data_ptr() if (nk and want_eval) else None)
    torch.cuda.synchronize()
    idx = b["idx"][:n, :k].cpu().numpy() if k else None
    cnt = b["cnt"][:n].cpu().numpy() if k else None
    nrm = b["nrm"][:n].cpu().numpy() if nk else None
    ev = b["ev"][:n].cpu().numpy() if (nk and want_eval) else None
    return idx, cnt, nrm, ev, plan


def single_cloud(cloud, k, nk, dmax=np.inf):
    """The plain single-cloud GPU path: one tree, self-query kNN of every
    valid row (as indices), normals of every row."""
    from pptk_amd import capi
    n = len(cloud)
    idx = np.full((n, max(k, 1)), -1, np.int32)[:, :k]
    cnt = np.zeros(n, np.int32)
    nrm = ev = None
    if n == 0:
        return idx, cnt, np.zeros((0, 3), cloud.dtype), np.zeros(0, cloud.dtype)
    tree = capi.Tree.build(cloud)
    valid = np.flatnonzero(np.all(np.isfinite(cloud), axis=1)).astype(np.int32)
    if k and len(valid):
        ii, cc = capi.knn(tree, valid, k, dmax)
        idx[valid] = ii
        cnt[valid] = cc
    if nk:
        nrm, ev, _, _ = capi.normals(tree, nk, dmax, rows=None, num_eigen=1)
    return idx, cnt, nrm, ev


def compare_with_single(clouds, idx, cnt, nrm, ev, k, nk, dmax=np.inf):
    off = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    for c, cloud in enumerate(clouds):
        a, b = off[c], off[c + 1]
        si, sc, sn, se = single_cloud(cloud, k, nk, dmax)
        if k:
            assert np.array_equal(cnt[a:b], sc), c
            assert np.array_equal(idx[a:b], si), c
        if nk:
            if not np.array_equal(nrm[a:b], sn, equal_nan=True):
                d = np.abs(nrm[a:b] - sn)
                raise AssertionError("cloud %d (n=%d): normals differ, max |d| = %g" %
                                     (c, b - a, np.nanmax(d)))
            assert np.array_equal(ev[a:b], se, equal_nan=True), c


def test_batch_f32_knn16_normals16_matches_single_cloud(gpu):
    rng = np.random.default_rng(11)
    clouds = _edge_clouds(rng)
    idx, cnt, nrm, ev, plan = run_batch(clouds, 3, np.float32, 16, 16)
    inf = plan.info()
    assert inf.num_clouds == len(clouds)
    assert inf.num_valid == sum(int(np.all(np.isfinite(c), axis=1).sum()) for c in clouds)
    compare_with_single(clouds, idx, cnt, nrm, ev, 16, 16)


def test_batch_against_oracle(gpu):
    """Sampled clouds against the oracle: kNN rows bit-exact, normals within
    0.5 degrees of FP64 (and the same neighbourhood sizes)."""
    rng = np.random.default_rng(12)
    clouds = _edge_clouds(rng)
    idx, cnt, nrm, ev, _ = run_batch(clouds, 3, np.float32, 16, 16)
    off = np.concatenate([[0], np.cumsum([len(c) for c in clouds])])
    checked = 0
    for c, cloud in enumerate(clouds):
        n = len(cloud)
        if n == 0 or not np.any(np.all(np.isfinite(cloud), axis=1)):
            continue
        a = off[c]
        ot = O.OracleTree(cloud)
        valid = np.flatnonzero(np.all(np.isfinite(cloud), axis=1)).astype(np.int32)
        wi, wc = ot.knn(valid, 16)
        assert np.array_equal(cnt[a + valid], wc), c
        assert np.array_equal(idx[a + valid], wi), c
        fev, few, fsz = ot.normals(16, np.inf)
        ok = fsz > 2
        ang = O.angle_deg(nrm[a:a + n][ok], fev[ok, 0, :])
        gap = (few[ok, 1] - few[ok, 0]) / np.where(few[ok, 2] > 0, few[ok, 2], 1)
        well = gap >= 1e-3
        if well.any():
            assert ang[well].max() < 0.5, (c, ang[well].max())
        checked += 1
    assert checked >= 40


def test_batch_graph_replay_and_new_input(gpu):
    """One plan, graph captured on the first run: a replay on new points
    gives the new results (nothing stale), and equals an eager plan."""
    rng = np.random.default_rng(13)
    clouds1 = _edge_clouds(rng)
    sizes = [len(c) for c in clouds1]
    clouds2 = [np.ascontiguousarray(rng.uniform(-5, 5, (s, 3)), np.float32) for s in sizes]
    r1 = run_batch(clouds1, 3, np.float32, 16, 16, use_graph=True)
    plan = r1[4]
    r2 = run_batch(clouds2, 3, np.float32, 16, 16, plan=plan)
    r1b = run_batch(clouds1, 3, np.float32, 16, 16, plan=plan)
    e2 = run_batch(clouds2, 3, np.float32, 16, 16, use_graph=False)
    for a, b in zip(r1[:4], r1b[:4]):
        assert np.array_equal(a, b, equal_nan=True)
    for a, b in zip(r2[:4], e2[:4]):
        assert np.array_equal(a, b, equal_nan=True)
    compare_with_single(clouds2, *r2[:4], 16, 16)


@pytest.mark.parametrize("k,nk,dmax", [(8, 16, np.inf), (32, 16, np.inf), (16, 0, np.inf),
                                       (0, 16, np.inf), (16, 16, 0.3), (4, 5, np.inf),
                                       (5, 32, np.inf)])
def test_batch_k_variants(gpu, k, nk, dmax):
    rng = np.random.default_rng(14 + k + nk)
    clouds = _edge_clouds(rng)
    dm = float(np.float32(dmax)) if np.isfinite(dmax) else dmax
    idx, cnt, nrm, ev, _ = run_batch(clouds, 3, np.float32, k, nk, dmax=dm)
    compare_with_single(clouds, idx, cnt, nrm, ev, k, nk, dm)


@pytest.mark.parametrize("leaf,qblock", [(8, 64), (32, 128), (64, 256), (1, 256)])
def test_batch_leaf_and_block_do_not_change_results(gpu, leaf, qblock):
    rng = np.random.default_rng(21)
    clouds = _edge_clouds(rng)
    base = run_batch(clouds, 3, np.float32, 16, 16)
    got = run_batch(clouds, 3, np.float32, 16, 16, leaf=leaf, qblock=qblock)
    for a, b in zip(base[:4], got[:4]):
        assert np.array_equal(a, b, equal_nan=True)


def test_batch_f64(gpu):
    rng = np.random.default_rng(15)
    clouds = [c + np.array([512000.0, 5400000.0, 300.0]) for c in
              _edge_clouds(rng, dtype=np.float64)]
    clouds = [np.ascontiguousarray(c) for c in clouds]
    idx, cnt, nrm, ev, _ = run_batch(clouds, 3, np.float64, 16, 16)
    compare_with_single(clouds, idx, cnt, nrm, ev, 16, 16)


@pytest.mark.parametrize("dim", [2, 4])
def test_batch_other_dims_knn(gpu, dim):
    rng = np.random.default_rng(16 + dim)
    clouds = _edge_clouds(rng, dim=dim)
    idx, cnt, _, _, _ = run_batch(clouds, dim, np.float32, 16, 0)
    compare_with_single(clouds, idx, cnt, None, None, 16, 0)


def test_batch_single_cloud_equals_tree(gpu):
    """A batch of one cloud is the plain path on that cloud."""
    rng = np.random.default_rng(17)
    cloud = np.ascontiguousarray(rng.uniform(0, 10, (200_000, 3)), np.float32)
    idx, cnt, nrm, ev, _ = run_batch([cloud], 3, np.float32, 16, 16)
    compare_with_single([cloud], idx, cnt, nrm, ev, 16, 16)


def test_batch_argument_errors(gpu):
    from pptk_amd import capi
    with pytest.raises(ValueError):
        capi.Batch([10, 20], k=33)
    with pytest.raises(ValueError):
        capi.Batch([10, 20], k=16, normals_k=33)
    with pytest.raises(ValueError):
        capi.Batch([10, 20], dim=2, normals_k=16)
    with pytest.raises(ValueError):
        capi.Batch([], k=16)
    with pytest.raises(ValueError):
        capi.Batch([10, -1], k=16)
    with pytest.raises(TypeError):
        capi.Batch([10], dtype=np.int32)
    with pytest.raises(ValueError):
        capi.Batch([10], query_block=96)
