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


def _hand_tree(pts, cuts):
    """Persisted float32 3-D tree written by hand: the points in x order, a
    root over all of them and a chain of leaves with the given tree-position
    cuts (leaves far larger than any bucket size the builder makes)."""
    hdr = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"),
                    ("dim", "<i4"), ("dtype", "<i4"), ("dist", "<i4"), ("leaf", "<i4"),
                    ("depth", "<i4"), ("pad", "<i4"),
                    ("n_input", "<i8"), ("n_valid", "<i8"), ("nn", "<i8"), ("nleaves", "<i8"),
                    ("nroots", "<i8"),
                    ("lo", "<f8", 4), ("hi", "<f8", 4), ("mlo", "<f8", 4), ("mscale", "<f8", 4),
                    ("off_coords", "<i8"), ("off_orig", "<i8"), ("off_fwd", "<i8"),
                    ("off_nodes", "<i8"), ("off_boxes", "<i8"), ("off_leafof", "<i8"),
                    ("total", "<i8")])
    n = len(pts)
    orig = np.argsort(pts[:, 0], kind="stable").astype(np.int32)
    tp = np.ascontiguousarray(pts[orig])
    fwd = np.empty(n, np.int32)
    fwd[orig] = np.arange(n, dtype=np.int32)
    # root -> (leaf [0, c0), inner [c0, n)) -> (leaf [c0, c1), inner [c1, n)) ...
    nodes = [[0, n, 1, -1]]
    bounds = [0] + list(cuts) + [n]
    parent = 0
    for i in range(len(bounds) - 2):
        a, b = bounds[i], bounds[i + 1]
        me = len(nodes)
        last = i == len(bounds) - 3
        nodes.append([a, b, -1, parent])
        nodes.append([b, n, -1 if last else me + 2, parent])
        nodes[parent][2] = me
        parent = me + 1
    nodes = np.array(nodes, np.int32)
    boxes = np.zeros((len(nodes), 2, 3), np.float32)
    leafof = np.zeros(n, np.int32)
    depth = np.zeros(len(nodes), np.int32)
    for i, (a, b, c, p) in enumerate(nodes):
        boxes[i, 0] = tp[a:b].min(0)
        boxes[i, 1] = tp[a:b].max(0)
        if c < 0:
            leafof[a:b] = i
        if p >= 0:
            depth[i] = depth[p] + 1
    h = np.zeros(1, hdr)
    h["magic"] = b"PPTKTREE"
    h["version"] = 1
    h["header_bytes"] = hdr.itemsize
    h["dim"], h["dtype"], h["dist"], h["leaf"], h["depth"] = 3, 0, 0, 256, depth.max() + 1
    h["n_input"], h["n_valid"], h["nn"] = n, n, len(nodes)
    h["nleaves"], h["nroots"] = int(np.sum(nodes[:, 2] < 0)), 1
    h["lo"][0, :3] = pts.min(0)
    h["hi"][0, :3] = pts.max(0)
    h["mlo"][0, :3] = pts.min(0)
    h["mscale"][0, :3] = 1.0
    o = (hdr.itemsize + 7) & ~7
    sections = [("off_coords", tp), ("off_orig", orig), ("off_fwd", fwd),
                ("off_nodes", nodes), ("off_boxes", boxes), ("off_leafof", leafof)]
    for name, arr in sections:
        h[name] = o
        o = (o + arr.nbytes + 7) & ~7
    h["total"] = o
    buf = np.zeros(o, np.uint8)
    buf[:hdr.itemsize] = np.frombuffer(h.tobytes(), np.uint8)
    for name, arr in sections:
        off = int(h[name][0])
        buf[off:off + arr.nbytes] = np.frombuffer(arr.tobytes(), np.uint8)
    return bytes(buf)


def test_restored_tree_with_leaves_above_256_points(gpu, monkeypatch):
    """Leaves of 700-1100 points and whole-range copies of up to 3000: the
    r-near scan and copy loops take such ranges 256 points at a time. Sets
    against the oracle, for point and index queries, replayed and walked
    fills; kNN rows too."""
    from pptk_amd import capi
    rng = np.random.default_rng(43)
    pts = rng.random((3000, 3)).astype(np.float32)
    r = capi.Tree.deserialize(_hand_tree(pts, [700, 1800]))
    assert r.info.num_nodes == 5
    ot = O.OracleTree(pts)
    qp = np.ascontiguousarray(rng.random((300, 3)).astype(np.float32))
    qi = rng.choice(len(pts), 300, replace=False).astype(np.int32)
    for chunks in (None, "0"):  # event lists kept (replay) / none (the fill walks)
        if chunks is None:
            monkeypatch.delenv("PPTK_RNEAR_EVENT_CHUNKS", raising=False)
        else:
            monkeypatch.setenv("PPTK_RNEAR_EVENT_CHUNKS", chunks)
        for dm in (0.15, 0.6, 2.0):
            dmax = float(np.float32(dm))
            for q in (qp, qi):
                off, idx = capi.rnear(r, q, dmax)
                woff, widx = ot.rnear(q, dmax)
                assert np.array_equal(off, woff), (chunks, dm, q.dtype)
                assert O.csr_first_mismatch(off, idx, woff, widx) == -1, (chunks, dm, q.dtype)
    gi, gc = capi.knn(r, None, 16)
    wi, wc = ot.knn(None, 16)
    assert np.array_equal(gi, wi) and np.array_equal(gc, wc)
