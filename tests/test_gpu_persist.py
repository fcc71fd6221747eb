"""GPU: persistent tree format round trips (pptk_tree_serialize /
pptk_tree_deserialize). A restored tree answers every query exactly as the
tree it was saved from, and re-serialising it gives the same bytes."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cases():
    from pptk_amd import scenes
    rng = np.random.default_rng(41)
    lid = scenes.lidar_scene(60_000, seed=41)
    f64 = rng.normal(size=(20_000, 3)) + np.array([512000.0, 5400000.0, 300.0])
    nan = rng.random((5000, 3)).astype(np.float32)
    nan[::17, 1] = np.nan
    return [("lidar_f32", lid), ("f64", f64), ("nan_rows", nan),
            ("dim2", rng.random((8000, 2)).astype(np.float32)),
            ("int16", rng.integers(-300, 300, (6000, 4)).astype(np.int16))]


@pytest.mark.parametrize("case", range(5))
def test_roundtrip_answers_identically(gpu, case, tmp_path):
    from pptk_amd import capi
    name, pts = _cases()[case]
    t = capi.Tree.build(pts)
    blob = t.serialize()
    path = tmp_path / (name + ".pptktree")
    t.save(str(path))
    assert path.read_bytes() == blob
    r = capi.Tree.load(str(path))
    assert r.serialize() == blob
    a, b = t.info, r.info
    assert (a.num_valid, a.num_nodes, a.leaf_size, a.dim) == (b.num_valid, b.num_nodes,
                                                              b.leaf_size, b.dim)
    k = 12
    assert all(np.array_equal(x, y) for x, y in zip(capi.knn(t, None, k), capi.knn(r, None, k)))
    rng = np.random.default_rng(case)
    q = pts[rng.choice(len(pts), 500)].astype(np.float64)
    q = q[np.all(np.isfinite(q), axis=1)]
    assert all(np.array_equal(x, y) for x, y in zip(capi.knn(t, q, k), capi.knn(r, q, k)))
    valid = np.flatnonzero(np.all(np.isfinite(pts), axis=1))[:2000]
    dm = float(np.float32(np.ptp(pts[valid].astype(np.float64), axis=0).max() * 0.02))
    o1, l1 = capi.rnear(t, valid, dm)
    o2, l2 = capi.rnear(r, valid, dm)
    assert np.array_equal(o1, o2) and np.array_equal(l1, l2)
    if pts.shape[1] == 3 and pts.dtype in (np.float32, np.float64):
        n1 = capi.normals(t, 16, np.inf)
        n2 = capi.normals(r, 16, np.inf)
        assert np.array_equal(n1[0], n2[0], equal_nan=True)
    # and the restored tree still matches the oracle
    wi, wc = O.OracleTree(pts).knn(None, k)
    gi, gc = capi.knn(r, None, k)
    assert np.array_equal(gi, wi) and np.array_equal(gc, wc)


def test_damaged_blob_is_rejected(gpu):
    from pptk_amd import capi
    rng = np.random.default_rng(42)
    t = capi.Tree.build(rng.random((3000, 3)).astype(np.float32))
    blob = bytearray(t.serialize())
    # break a child link in the node section: header says where it starts
    off_nodes = int(np.frombuffer(bytes(blob[232:240]), np.int64)[0])
    nodes = np.frombuffer(bytes(blob[off_nodes:off_nodes + 16 * 4]), np.int32).copy()
    nodes[2] = 2  # root's child -> an even id
    blob[off_nodes:off_nodes + 16 * 4] = nodes.tobytes()
    with pytest.raises(ValueError):
        capi.Tree.deserialize(bytes(blob))
