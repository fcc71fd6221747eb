"""GPU against the committed golden vectors: pptk's own k-d tree outputs
(tests/golden/ref_*.npz) and the reference tests' known answers."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _mat(lst, k):
    out = np.full((len(lst), k), -1, np.int32)
    for i, x in enumerate(lst):
        out[i, :len(x)] = x
    return out, np.array([len(x) for x in lst], np.int32)


def test_known_answer_knn(gpu):
    """pptk tests/test_points.py:299-308 through kdtree._query."""
    import pptk_amd
    g = np.load(os.path.join(GOLD, "known_answer_knn.npz"))
    tree = pptk_amd.kdtree._build(g["P"])
    R = np.vstack(pptk_amd.kdtree._query(tree, g["Q"], k=10))
    assert np.array_equal(R, g["J"])


@pytest.mark.parametrize("name", ["ref_uniform_f64.npz", "ref_lidar_f32.npz"])
def test_reference_vectors(gpu, name):
    import pptk_amd
    g = np.load(os.path.join(GOLD, name))
    kd = pptk_amd.kdtree
    tree = kd._build(g["points"])
    k = int(g["k_self"])
    idx, cnt = _mat(kd._query(tree, None, k=k), k)
    assert np.array_equal(cnt, g["self_cnt"]) and np.array_equal(idx, g["self_idx"])
    kq = int(g["k_q"])
    idx, cnt = _mat(kd._query(tree, g["queries"], k=kq, dmax=float(g["dmax_q"])), kq)
    assert np.array_equal(cnt, g["q_cnt"]) and np.array_equal(idx, g["q_idx"])
    res = kd._query(tree, None, k=None, dmax=float(g["r_near"]))
    off, ridx = g["rnear_off"], g["rnear_idx"]
    assert len(res) == len(off) - 1
    for i, x in enumerate(res):
        assert np.array_equal(np.sort(x), ridx[off[i]:off[i + 1]]), i


def test_collinear_normals_fixture(gpu):
    import pptk_amd
    g = np.load(os.path.join(GOLD, "collinear_normals.npz"))
    ev, ew, sz = pptk_amd.estimate_normals(g["points"], int(g["k"]), float(g["r"]),
                                           output_eigenvalues=True, output_all_eigenvectors=True,
                                           output_neighborhood_sizes=True, verbose=False)
    assert np.array_equal(sz, g["sizes"])
    assert np.allclose(ew[:, -1], g["evals_max"])
    assert np.allclose(ev[:, -1, :], g["evecs_max"])
