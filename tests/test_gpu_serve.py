"""GPU: the serving loop (pptk_amd.serve) on the device.

Every request's result must be identical alone (one slot, nothing else in
flight) and in a mix (eight slots, requests overlapping on the device, more
traffic on other clouds, trees resident across follow-ups); a sample of the
results is also checked against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _clouds():
    from pptk_amd import scenes
    sizes = [12_000, 30_000, 55_000, 90_000, 140_000, 200_000]
    return [scenes.lidar_scene(n, seed=100 + i) for i, n in enumerate(sizes)]


def _requests(clouds, rng):
    from pptk_amd import serve
    reqs = []
    kinds = [("knn", 8, np.inf), ("knn", 16, np.inf), ("knn", 32, np.inf),
             ("rnear", 16, 0.05), ("rnear", 16, 0.2), ("normals", 16, np.inf)]
    for i in range(24):
        c = i % len(clouds)
        kind, k, dmax = kinds[i % len(kinds)]
        q = None
        if kind == "rnear":
            q = np.sort(rng.choice(len(clouds[c]), 4000, replace=False)).astype(np.int32)
        dm = float(np.float32(dmax)) if np.isfinite(dmax) else dmax
        # the first request of a cloud brings its points; later ones are
        # follow-ups on the resident tree
        pts = clouds[c] if i < len(clouds) else None
        reqs.append(serve.Request(kind, "c%d" % c, pts, k=k, dmax=dm, queries=q, tag=i))
    return reqs


def _np(x):
    if isinstance(x, tuple):
        return tuple(_np(y) for y in x)
    return x.cpu().numpy() if hasattr(x, "cpu") else x


def _same(a, b):
    a, b = _np(a), _np(b)
    if isinstance(a, tuple):
        return len(a) == len(b) and all(np.array_equal(x, y) for x, y in zip(a, b))
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("outputs", ["host", "device"])
def test_results_identical_alone_and_in_mix(gpu, outputs):
    from pptk_amd import serve
    rng = np.random.default_rng(31)
    clouds = _clouds()
    reqs = _requests(clouds, rng)
    s1 = serve.Server(serve.GpuBackend(1, outputs=outputs), slots=1)
    alone = []
    for r in reqs:
        alone.append(_np(s1.submit(r).result(timeout=120)))
    s1.close()

    s8 = serve.Server(serve.GpuBackend(8, outputs=outputs), slots=8)
    # background traffic on other clouds, interleaved with the requests
    extra = [serve.Request("knn", "x%d" % i, c, k=16) for i, c in enumerate(clouds)]
    extra += [serve.Request("normals", "x%d" % i, None, k=16) for i in range(len(clouds))]
    order = []
    for i, r in enumerate(reqs):
        order.append((r, True))
        if i < len(extra):
            order.append((extra[i], False))
    tickets = [(s8.submit(r), mine) for r, mine in order]
    mixed = [t.result(timeout=120) for t, mine in tickets if mine]
    for t, mine in tickets:
        t.result(timeout=120)
    assert s8.memory()["resident_trees"] >= len(clouds)
    s8.close()
    for r, a, b in zip(reqs, alone, mixed):
        assert _same(a, b), (r.tag, r.kind)


def test_served_results_match_oracle(gpu):
    from pptk_amd import serve
    rng = np.random.default_rng(32)
    clouds = _clouds()[:3]
    s = serve.Server(serve.GpuBackend(4), slots=4)
    t_knn = s.submit(serve.Request("knn", "a", clouds[1], k=16))
    q = np.sort(rng.choice(len(clouds[2]), 3000, replace=False)).astype(np.int32)
    dm = float(np.float32(0.2))
    t_rn = s.submit(serve.Request("rnear", "b", clouds[2], dmax=dm, queries=q))
    t_nm = s.submit(serve.Request("normals", "c", clouds[0], k=16))
    idx, cnt = t_knn.result(timeout=120)
    off, lst = t_rn.result(timeout=120)
    nrm = t_nm.result(timeout=120)
    s.close()
    wi, wc = O.OracleTree(clouds[1]).knn(None, 16)
    assert np.array_equal(idx, wi) and np.array_equal(cnt, wc)
    woff, wl = O.OracleTree(clouds[2]).rnear(q, dm)
    assert np.array_equal(off, woff)
    for i in range(len(q)):
        assert np.array_equal(np.sort(lst[off[i]:off[i + 1]]), wl[woff[i]:woff[i + 1]])
    fev, few, fsz = O.OracleTree(clouds[0]).normals(16, np.inf)
    gap = (few[:, 1] - few[:, 0]) / np.maximum(few[:, 2], 1e-300)
    ang = O.angle_deg(nrm, fev[:, 0, :])
    assert ang[gap >= 1e-3].max() < 0.5


def test_checkpoint_restore_on_device(gpu, tmp_path):
    from pptk_amd import serve
    clouds = _clouds()[:3]
    s = serve.Server(serve.GpuBackend(2), slots=2)
    first = [s.submit(serve.Request("knn", "c%d" % i, c, k=16)).result(timeout=120)
             for i, c in enumerate(clouds)]
    assert s.checkpoint(str(tmp_path)) == 3
    s.close()
    s2 = serve.Server(serve.GpuBackend(2), slots=2)
    assert s2.restore(str(tmp_path)) == 3
    again = [s2.submit(serve.Request("knn", "c%d" % i, None, k=16)).result(timeout=120)
             for i in range(3)]
    assert not any(t.built for t in s2.completed)
    s2.close()
    for a, b in zip(first, again):
        assert _same(a, b)
