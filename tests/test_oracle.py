"""CPU: pins the oracle before it is trusted as the GPU checker.

The oracle (oracle/pptk_oracle.cpp) is checked against
  * the reference tests' known answers (pptk tests/test_points.py:299-308,
    tests/test_estimate_normals.py:38-183) via tests/golden fixtures;
  * pptk's own k-d tree compiled from the reference sources (golden fixtures
    written by tests/golden/make_golden.py, and oracle/_ref live when built);
  * exhaustive NumPy search evaluating the reference distance formula;
  * LAPACK (numpy.linalg.eigh) for the FP64 normals;
so that a dropped term, a transposed operand, a wrong tie order or a wrong
comparison (< vs <=) fails at least one check.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


# -------------------------------------------------------- distance formula --

def test_pair_dist2_is_left_to_right_without_fma():
    # In float32, (a + b) + c differs from a + (b + c) for these terms.
    q = np.zeros((1, 3), np.float32)
    p = np.array([[1.0, 2.0 ** -12, 2.0 ** -12]], np.float32)
    d = O.pair_dist2(q, p, np.float32)[0]
    left = (np.float32(1.0) + np.float32(2.0 ** -24)) + np.float32(2.0 ** -24)
    assert d == left == np.float32(1.0)
    right = np.float32(1.0) + (np.float32(2.0 ** -24) + np.float32(2.0 ** -24))
    assert right != left  # the order matters, and the oracle uses the left one


def test_known_answer_knn():
    """pptk tests/test_points.py:299-308."""
    g = load("known_answer_knn.npz")
    idx, cnt = O.OracleTree(g["P"]).knn(g["Q"], 10)
    assert np.array_equal(idx, g["J"])
    b, _ = O.brute_knn(g["P"], g["Q"], 10)
    assert np.array_equal(b, g["J"])


@pytest.mark.parametrize("name", ["ref_uniform_f64.npz", "ref_lidar_f32.npz"])
def test_oracle_matches_reference_fixtures(name):
    g = load(name)
    t = O.OracleTree(g["points"])
    idx, cnt = t.knn(None, int(g["k_self"]))
    assert np.array_equal(cnt, g["self_cnt"])
    assert np.array_equal(idx, g["self_idx"])
    idx, cnt = t.knn(g["queries"], int(g["k_q"]), dmax=float(g["dmax_q"]))
    assert np.array_equal(cnt, g["q_cnt"])
    assert np.array_equal(idx, g["q_idx"])
    off, ridx = t.rnear(None, float(g["r_near"]))
    assert np.array_equal(off, g["rnear_off"])
    assert np.array_equal(ridx, g["rnear_idx"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_matches_reference_live(dtype):
    rng = np.random.default_rng(42)
    pts = rng.normal(size=(20_000, 3)).astype(dtype)
    q = rng.normal(size=(2000, 3)).astype(dtype)
    t, r = O.OracleTree(pts), O.RefTree(pts)
    assert np.array_equal(t.knn(None, 12)[0], r.knn(None, 12)[0])
    assert np.array_equal(t.knn(q, 7, dmax=0.3)[0], r.knn(q, 7, dmax=0.3)[0])
    a, b = t.rnear(q, 0.2), r.rnear(q, 0.2)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ------------------------------------------------------------ brute force --

@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16])
@pytest.mark.parametrize("dim", [2, 3, 4])
def test_oracle_vs_brute_knn(dtype, dim):
    rng = np.random.default_rng(dim)
    pts = (rng.random((3000, dim)) * 100).astype(dtype)
    t = O.OracleTree(pts)
    for k in (1, 4, 9):
        idx, cnt = t.knn(None, k)
        b, bc = O.brute_knn(pts, pts, k, self_rows=np.arange(3000))
        assert np.array_equal(idx, b) and np.array_equal(cnt, bc)
    q = (rng.random((200, dim)) * 100).astype(np.float64)
    idx, cnt = t.knn(q, 6, dmax=7.0)
    b, bc = O.brute_knn(pts, q, 6, dmax=7.0)
    assert np.array_equal(idx, b) and np.array_equal(cnt, bc)


def test_oracle_ties_prefer_smaller_index():
    g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8)), -1).reshape(-1, 3)
    pts = np.concatenate([g, g[::5]]).astype(np.float32)
    t = O.OracleTree(pts)
    for k in (1, 6, 7, 19, 30):
        idx, cnt = t.knn(None, k)
        b, bc = O.brute_knn(pts, pts, k, self_rows=np.arange(len(pts)))
        assert np.array_equal(idx, b), k
    # a point exactly at the k-th distance but with a smaller index wins
    pts2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0]], np.float64)
    idx, _ = O.OracleTree(pts2).knn(np.zeros((1, 3)), 2)
    assert list(idx[0]) == [0, 1]
    idx, _ = O.OracleTree(pts2).knn(np.array([0]), 2)
    assert list(idx[0]) == [1, 2]


def test_oracle_strict_radius():
    pts = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float64)
    t = O.OracleTree(pts)
    idx, cnt = t.knn(np.zeros((1, 3)), 3, dmax=1.0)
    assert cnt[0] == 1  # d = 1 is not < 1
    off, ridx = t.rnear(np.zeros((1, 3)), 2.0)
    assert list(ridx) == [0, 1]


def test_oracle_nan_rows_and_queries():
    rng = np.random.default_rng(3)
    pts = rng.random((2000, 3))
    pts[::17, 1] = np.nan
    pts[5] = np.inf
    t = O.OracleTree(pts)
    valid = np.all(np.isfinite(pts), 1)
    assert t.num_valid == valid.sum()
    idx, cnt = t.knn(np.arange(2000), 5)
    assert np.all(cnt[~valid] == 0)
    P = np.where(valid[:, None], pts, np.nan)
    b, bc = O.brute_knn(P, np.nan_to_num(pts, nan=np.nan), 5, self_rows=np.arange(2000))
    assert np.array_equal(idx[valid], b[valid])
    q = np.array([[np.nan, 0, 0], [0.5, 0.5, 0.5]])
    idx, cnt = t.knn(q, 3)
    assert cnt[0] == 0 and cnt[1] == 3


def test_oracle_vs_brute_rnear():
    rng = np.random.default_rng(5)
    pts = rng.random((3000, 3)).astype(np.float32)
    t = O.OracleTree(pts)
    off, idx = t.rnear(None, 0.08)
    br = O.brute_rnear(pts, pts, 0.08, self_rows=np.arange(3000))
    for i in range(3000):
        assert np.array_equal(idx[off[i]:off[i + 1]], br[i])
    q = rng.random((300, 3))
    off, idx = t.rnear(q, 0.1)
    br = O.brute_rnear(pts, q, 0.1)
    for i in range(300):
        assert np.array_equal(idx[off[i]:off[i + 1]], br[i])


# --------------------------------------------------------------- normals --

def test_oracle_normals_collinear_closed_form():
    g = load("collinear_normals.npz")
    ev, ew, sz = O.OracleTree(g["points"]).normals(int(g["k"]), float(g["r"]))
    assert np.array_equal(sz, g["sizes"])
    assert np.allclose(ew[:, -1], g["evals_max"])
    assert np.allclose(ev[:, -1, :], g["evecs_max"])


def test_oracle_normals_vs_lapack():
    rng = np.random.default_rng(7)
    pts = rng.normal(size=(5000, 3)) * [3.0, 1.0, 0.2]
    t = O.OracleTree(pts)
    ev, ew, sz = t.normals(16, np.inf)
    idx, cnt = t.knn(pts, 16)
    lev, lew = O.normals_fp64(pts, idx, cnt)
    assert np.array_equal(sz, cnt)
    assert np.allclose(ew, lew, rtol=1e-10, atol=1e-14)
    gap = (lew[:, 1] - lew[:, 0]) / lew[:, 2]
    ok = gap > 1e-6
    for j in range(3):
        assert O.angle_deg(ev[ok, j], lev[ok, j]).max() < 1e-4
    # orthonormal eigenvector basis, largest component positive
    G = np.einsum("mij,mkj->mik", ev, ev)
    assert np.allclose(G, np.eye(3), atol=1e-12)
    big = np.take_along_axis(ev, np.abs(ev).argmax(-1)[..., None], -1)
    assert np.all(big > 0)


def test_oracle_normals_radius_modes():
    rng = np.random.default_rng(8)
    pts = rng.random((3000, 3))
    t = O.OracleTree(pts)
    ev, ew, sz = t.normals(-1, 0.1)
    off, idx = t.rnear(pts, 0.1)
    assert np.array_equal(sz, np.diff(off))
    ev, ew, sz = t.normals(8, 0.05)
    i2, c2 = t.knn(pts, 8, dmax=0.05)
    assert np.array_equal(sz, c2)
    lev, lew = O.normals_fp64(pts, i2, c2)
    ok = c2 >= 3
    assert np.allclose(ew[ok], lew[ok], rtol=1e-9, atol=1e-15)


def test_angle_helper():
    a = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    b = np.array([[-1.0, 0, 0], [0, np.cos(0.01), np.sin(0.01)]])
    ang = O.angle_deg(a, b)
    assert ang[0] == 0 and np.isclose(ang[1], np.degrees(0.01))
