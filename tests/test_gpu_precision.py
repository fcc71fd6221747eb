"""GPU: the mixed-precision normals' FP64 fallback lane.

Isolated float32 clusters of 16 points (pptk_amd.scenes.gap_clusters) have
covariances with eigenvalues (1, 1 + 4g, 4) in random orientations, so each
point's 16-neighbourhood (k=16, or the r=10 ball) is its cluster and its
eigengap (l1 - l0) / l2 is the chosen g. Gaps 0 .. 2e-4 lie below the 1e-3
threshold of PPTK_PREC_MIXED, gaps 3e-3 .. 0.5 above it, both far enough
from it that float32 rounding (~1e-7) cannot move a cluster across.

* The fallback count equals the number of points in below-threshold
  clusters, exactly.
* Those rows equal the PPTK_PREC_FP64 rows bit for bit (the fallback
  recomputes the covariance in FP64 from the same neighbours), and lie within
  1e-5 degrees of LAPACK's eigenvector: an FP32 solve would be off by about
  1e-7 / g radians, 0.6 degrees at g = 1e-5.
* Rows above the threshold stay within the 0.5 degree tolerance.
Reference: estimate_normals.cpp:84-89 (centred covariance, Eigen's
SelfAdjointEigenSolver), here pinned by numpy.linalg.eigh in FP64.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GAPS = [0.0, 1e-6, 1e-5, 1e-4, 2e-4, 3e-3, 1e-2, 0.1, 0.5]
PER_GAP = 24


def _clusters(seed):
    from pptk_amd import scenes
    return scenes.gap_clusters(GAPS, per_gap=PER_GAP, seed=seed)


def _fp64_reference(pts):
    ot = O.OracleTree(pts)
    nb, cnt = ot.knn(pts, 16)
    assert np.all(cnt == 16)
    return O.normals_fp64(pts, nb, cnt)


@pytest.mark.parametrize("rows_mode", ["all", "list"])
@pytest.mark.parametrize("k,r", [(16, np.inf), (-1, 10.0)])
def test_fallback_rows_equal_fp64_mode(gpu, rows_mode, k, r):
    from pptk_amd import capi
    pts, g = _clusters(seed=1)
    n = len(pts)
    low = g < 1e-3
    tree = capi.Tree.build(pts)
    rows = None if rows_mode == "all" else np.arange(n, dtype=np.int32)
    mix = capi.normals(tree, k, r, rows=rows, num_eigen=3, precision=capi.PREC_MIXED)
    f64 = capi.normals(tree, k, r, rows=rows, num_eigen=3, precision=capi.PREC_FP64)
    assert np.all(mix[2] == 16) and np.all(f64[2] == 16)
    # every below-threshold row falls back; no other row does
    assert mix[3] == int(low.sum())
    assert f64[3] == 0
    assert np.array_equal(mix[0][low], f64[0][low])
    assert np.array_equal(mix[1][low], f64[1][low])
    fev, few = _fp64_reference(pts)
    ang = O.angle_deg(mix[0][:, 0], fev[:, 0])
    assert ang[low].max() < 1e-5, ang[low].max()
    assert ang[~low].max() < 0.5, ang[~low].max()
    # the eigenvalues of fallback rows are FP64 values rounded once to the
    # output dtype (float32): within one float32 ulp of LAPACK's
    assert np.allclose(mix[1][low], few[low], rtol=1.2e-7, atol=0)
    # the largest eigenvector is well defined in every row
    big = O.angle_deg(mix[0][:, 2], fev[:, 2])
    assert big.max() < 1e-3
    tree.free()


def test_fallback_count_normals_only_and_estimate_normals(gpu):
    """num_eigen=1 (the normal only) uses the same rule through the drop-in."""
    import pptk_amd
    from pptk_amd import capi
    pts, g = _clusters(seed=2)
    low = g < 1e-3
    tree = capi.Tree.build(pts)
    ev, ew, sz, nfb = capi.normals(tree, 16, np.inf, num_eigen=1, precision=capi.PREC_MIXED)
    assert nfb == int(low.sum())
    fev, _ = _fp64_reference(pts)
    ang = O.angle_deg(ev, fev[:, 0])
    assert ang[low].max() < 1e-5 and ang[~low].max() < 0.5
    nrm = pptk_amd.estimate_normals(pts, 16, np.inf, verbose=False)
    assert np.array_equal(nrm, ev)
    tree.free()


def test_batched_fallback_matches_single_cloud(gpu):
    """The fused batch kernel takes the same fallback: its normals equal the
    single-cloud mixed normals bit for bit, clouds of clusters included."""
    import torch

    from pptk_amd import capi
    clouds = [_clusters(seed=s)[0] for s in (3, 4, 5)]
    sizes = [len(c) for c in clouds]
    want, nfb_want = [], 0
    for c in clouds:
        t = capi.Tree.build(c)
        ev, _, _, nfb = capi.normals(t, 16, np.inf, num_eigen=1, precision=capi.PREC_MIXED)
        want.append(ev)
        nfb_want += nfb
        t.free()
    pts_d = torch.from_numpy(np.concatenate(clouds)).cuda()
    N = pts_d.shape[0]
    idx_d = torch.empty((N, 16), dtype=torch.int32, device="cuda")
    nrm_d = torch.empty((N, 3), dtype=torch.float32, device="cuda")
    plan = capi.Batch(sizes, k=16, normals_k=16, precision=capi.PREC_MIXED, use_graph=False)
    torch.cuda.synchronize()
    plan.run(pts_d.data_ptr(), idx_d.data_ptr(), None, nrm_d.data_ptr(), None)
    torch.cuda.synchronize()
    assert plan.info().num_fp64 == nfb_want
    assert np.array_equal(nrm_d.cpu().numpy(), np.concatenate(want))
    plan.free()
