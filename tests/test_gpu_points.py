"""GPU: Points (pptk_amd.points), the host-side mirror of pptk's direct
caller of the k-d tree, on the device: nbhds/NBHDS against the oracle,
before and after modifications."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rows(lst, k):
    out = np.full((len(lst), k), -1, np.int32)
    for i, x in enumerate(lst):
        out[i, :len(x)] = x
    return out


def test_points_nbhds_and_rebuild_after_modification(gpu):
    from pptk_amd import points as P
    from pptk_amd import scenes
    base = scenes.lidar_scene(30_000, seed=51)
    x = P.points(base)
    got = _rows(x.nbhds(k=8), 8)
    want, _ = O.OracleTree(base).knn(None, 8)
    assert np.array_equal(got, want)
    assert not x._update_kd_tree()           # unchanged: the resident tree stays
    x[100:200] += np.float32(0.5)            # modify a block of points
    mod = np.array(base)
    mod[100:200] += np.float32(0.5)
    assert np.array_equal(np.asarray(x), mod)
    got = _rows(x.nbhds(k=8), 8)
    want, _ = O.OracleTree(mod).knn(None, 8)
    assert np.array_equal(got, want)
    q = np.arange(0, 30_000, 7)
    got = _rows(x.nbhds(queries=q, k=5, r=0.3), 5)
    want, _ = O.OracleTree(mod).knn(q, 5, dmax=float(np.float32(0.3)))
    assert np.array_equal(got, want)


def test_points_NBHDS_lazy_matches_direct(gpu):
    from pptk_amd import points as P
    rng = np.random.default_rng(52)
    x = P.points(rng.random((12_000, 3)))
    e = x.NBHDS(k=6, r=0.05)
    direct = x.nbhds(k=6, r=0.05)
    ev = e.evaluate()
    it = list(e)
    assert len(ev) == len(it) == len(direct) == 12_000
    for a, b, c in zip(ev, it, direct):
        assert np.array_equal(a, c) and np.array_equal(b, c)
    assert np.array_equal(e[11_999], direct[11_999])
