"""GPU: pptk_knn_normals, the self-query kNN rows and the PCA normals of one
tree from a single search per point, against the two calls whose outputs it
defines: pptk_knn with queries=None (kdtree._query(tree, k=k, dmax=dmax),
kdtree_wrapper.cpp:281-304) and pptk_normals with rows=None
(estimate_normals(P, nk, r), estimate_normals.cpp:64-111). Every output is
compared bit for bit (the same neighbour sets, summed in the same order),
and one case is also checked against the oracle's kNN and FP64 normals."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _separate(tree, k, nk, dmax, r, ne, prec):
    from pptk_amd import capi
    idx, cnt, dist = capi.knn(tree, None, k, dmax, with_dist=True)
    ev, ew, sz, nf = capi.normals(tree, nk, r, rows=None, num_eigen=ne, precision=prec)
    return idx, cnt, dist, ev, ew, sz, nf


def _fused(tree, k, nk, dmax, r, ne, prec):
    from pptk_amd import capi
    return capi.knn_normals(tree, k, nk, dmax=dmax, r=r, num_eigen=ne, precision=prec,
                            with_dist=True)


def _assert_same(a, b, what):
    names = ("idx", "cnt", "dist", "evecs", "evals", "sizes", "num_fp64")
    for n, x, y in zip(names, a, b):
        if isinstance(x, np.ndarray):
            assert x.shape == y.shape and x.dtype == y.dtype, (what, n)
            assert np.array_equal(x, y, equal_nan=x.dtype.kind == "f"), (what, n)
        else:
            assert x == y, (what, n)


def _scene(n, seed):
    from pptk_amd import scenes
    return scenes.lidar_scene(n, seed=seed)


CASES = [
    # k, nk, dmax, r, num_eigen, precision
    (16, 16, np.inf, np.inf, 1, 0),
    (8, 16, np.inf, np.inf, 1, 1),
    (16, 8, np.inf, np.inf, 3, 2),
    (5, 12, 0.05, 0.08, 1, 1),
    (12, 5, 0.08, 0.05, 3, 1),
    (32, 32, np.inf, np.inf, 1, 1),
    (3, 1, np.inf, np.inf, 1, 1),
    (1, 4, np.inf, 0.02, 1, 2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "k%d_nk%d_d%g_r%g_ne%d_p%d" % c)
def test_fused_equals_separate_calls(gpu, case):
    from pptk_amd import capi
    k, nk, dmax, r, ne, prec = case
    pts = _scene(300_000, 3)
    tree = capi.Tree.build(pts)
    _assert_same(_fused(tree, *case), _separate(tree, *case), case)


def test_fused_nan_rows_duplicates_and_zero_distance_ties(gpu):
    """Rows of NaN/inf points (kNN rows pre-filled empty, normals NaN), 24
    copies of one point and 30 distinct points whose squared distances
    underflow to 0 in float32: more than nk points tie with a point at d = 0
    ahead of it, so the point itself is not in its own neighbourhood."""
    from pptk_amd import capi
    rng = np.random.default_rng(11)
    base = rng.uniform(0, 1, (20_000, 3)).astype(np.float32)
    dup = np.repeat(base[:1], 24, axis=0)
    tiny = (np.arange(30, dtype=np.float32)[:, None] * np.float32(1e-25)
            * np.ones((1, 3), np.float32)).astype(np.float32)
    assert np.float32(1e-25) ** 2 == 0
    pts = np.concatenate([dup, tiny, base]).astype(np.float32)
    pts[rng.choice(len(pts), 300, replace=False), rng.integers(0, 3, 300)] = np.nan
    pts[7] = np.inf
    pts = pts[rng.permutation(len(pts))]
    tree = capi.Tree.build(pts)
    assert tree.info.num_valid < len(pts)
    for case in [(16, 16, np.inf, np.inf, 1, 1), (4, 16, np.inf, np.inf, 3, 2),
                 (20, 17, np.inf, np.inf, 1, 1)]:
        _assert_same(_fused(tree, *case), _separate(tree, *case), case)


def test_fused_float64_and_large_k_take_the_separate_searches(gpu):
    from pptk_amd import capi
    pts = _scene(50_000, 4)
    t64 = capi.Tree.build(pts.astype(np.float64))
    case = (16, 16, np.inf, np.inf, 1, 0)
    _assert_same(_fused(t64, *case), _separate(t64, *case), case)
    t32 = capi.Tree.build(pts)
    case = (40, 16, np.inf, np.inf, 1, 1)
    _assert_same(_fused(t32, *case), _separate(t32, *case), case)


def test_fused_against_the_oracle(gpu):
    """kNN rows bit-exact against the oracle's self query; normals within
    0.5 degrees of LAPACK on the oracle's 16-neighbourhoods (including the
    point itself)."""
    from pptk_amd import capi
    pts = _scene(60_000, 5)
    tree = capi.Tree.build(pts)
    idx, cnt, _, ev, ew, sz, _ = capi.knn_normals(tree, 16, 16)
    ot = O.OracleTree(pts)
    want, wcnt = ot.knn(None, 16)
    assert np.array_equal(cnt, wcnt)
    assert np.array_equal(idx, want)
    nb, nc = ot.knn(pts, 16)
    assert np.array_equal(sz, nc)
    fev, _ = O.normals_fp64(pts, nb, nc)
    ang = O.angle_deg(ev, fev[:, 0])
    assert float(ang.max()) < 0.5


def test_fused_argument_errors(gpu):
    from pptk_amd import capi
    tree = capi.Tree.build(_scene(2_000, 6))
    for kw in [dict(k=0, nk=16), dict(k=16, nk=0), dict(k=16, nk=16, r=0.0),
               dict(k=16, nk=16, dmax=-1.0), dict(k=16, nk=16, num_eigen=2)]:
        with pytest.raises(Exception):
            capi.knn_normals(tree, **kw)
    t2 = capi.Tree.build(np.random.default_rng(0).uniform(0, 1, (500, 2)).astype(np.float32))
    with pytest.raises(Exception):
        capi.knn_normals(t2, 4, 4)
