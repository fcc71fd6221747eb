"""Seeded synthetic point clouds for the benchmark workloads and tests.

* ``uniform_cube``: n float64 points i.i.d. uniform in [0,1)^3 with
  ``np.random.seed(seed)`` + ``np.random.rand`` (the reference tests' inputs,
  pptk/tests/test_points.py:283-285).
* ``lidar_scene``: a lidar-like scan with the sensor at the origin: ground
  plane z=0 over [-100,100]^2 m, 40 vertical wall rectangles (axis-aligned,
  length U[10,50] m, height U[3,15] m) and 200 vertical cylinders (radius
  U[0.2,1] m, height U[2,10] m). Points are drawn with surface density
  proportional to 1/max(range, 1 m)^2, the point budget split between
  surfaces in proportion to their integrated density, plus isotropic
  N(0, 0.01 m) noise.
* ``precision_stress``: the lidar scene in float64 translated by
  (512000, 5400000, 300) m plus a near-isotropic cluster of 2M points around
  c = (10, 10, 1) m (before translation) with standard deviation 0.05 m per
  axis. i.i.d. Gaussian points would give k=16 neighbourhoods with random
  covariances, almost never with a small eigengap; the cluster is therefore
  built from 125,000 groups of 16 points, each the isotropic 16-point design
  (``isotropic_design16``) in a random orientation, 0.4 mm across, jittered
  by a log-uniform 1e-9..1e-6 m per group, on the lattice points nearest c
  (3.6 mm apart): every neighbourhood is one group and its eigengap
  (l1 - l0) / l2 spreads from ~1e-6 to ~1e-2.
* ``gap_clusters``: isolated float32 16-point clusters whose covariance has
  eigenvalues (1, 1 + 4g, 4) in a random orientation, for chosen gaps g.
* ``batched_clouds``: clouds of log-uniform size, each a random mix of noisy
  planes and spheres in a 10 m cube.
"""
import numpy as np

R_MIN = 1.0


def uniform_cube(n=100_000, seed=0):
    np.random.seed(seed)
    return np.random.rand(n, 3)


def _weight(p):
    r = np.sqrt(np.sum(p * p, axis=1))
    return 1.0 / np.maximum(r, R_MIN) ** 2


def _wall_points(rng, axis, length, h, cx, cy, m):
    u = rng.random(m) - 0.5
    z = rng.random(m) * h
    if axis == 0:
        return np.stack([cx + u * length, np.full(m, cy), z], axis=1)
    return np.stack([np.full(m, cx), cy + u * length, z], axis=1)


def _cyl_points(rng, r, h, cx, cy, m):
    th = rng.random(m) * 2 * np.pi
    z = rng.random(m) * h
    return np.stack([cx + r * np.cos(th), cy + r * np.sin(th), z], axis=1)


def _sample_weighted(rng, gen, m, wmax):
    """m points from gen(count) thinned by _weight / wmax (rejection)."""
    out, have = [], 0
    while have < m:
        want = int((m - have) * 1.5) + 256
        p = gen(want)
        p = p[rng.random(want) < _weight(p) / wmax]
        out.append(p)
        have += len(p)
    return np.concatenate(out)[:m]


def lidar_scene(n=10_000_000, seed=1, noise=0.01, dtype=np.float32):
    rng = np.random.default_rng(seed)
    half = 100.0
    nw, nc = 40, 200
    w_axis = rng.integers(0, 2, nw)
    w_len = rng.uniform(10, 50, nw)
    w_h = rng.uniform(3, 15, nw)
    w_cx = rng.uniform(-90, 90, nw)
    w_cy = rng.uniform(-90, 90, nw)
    c_r = rng.uniform(0.2, 1.0, nc)
    c_h = rng.uniform(2, 10, nc)
    c_cx = rng.uniform(-95, 95, nc)
    c_cy = rng.uniform(-95, 95, nc)

    # Each surface: a sampler, its closest distance to the sensor (bound on
    # the weight) and its integrated density (area x mean weight, by MC).
    surf = []
    k = 4096
    for i in range(nw):
        along = w_cx[i] if w_axis[i] == 0 else w_cy[i]
        across = w_cy[i] if w_axis[i] == 0 else w_cx[i]
        gap = max(0.0, abs(along) - w_len[i] / 2)
        dmin = np.hypot(gap, across)
        gen = (lambda a, L, hh, x, y: (lambda m: _wall_points(rng, a, L, hh, x, y, m)))(
            w_axis[i], w_len[i], w_h[i], w_cx[i], w_cy[i])
        surf.append((gen, dmin, w_len[i] * w_h[i] * np.mean(_weight(gen(k)))))
    for i in range(nc):
        dmin = max(0.0, np.hypot(c_cx[i], c_cy[i]) - c_r[i])
        gen = (lambda rr, hh, x, y: (lambda m: _cyl_points(rng, rr, hh, x, y, m)))(
            c_r[i], c_h[i], c_cx[i], c_cy[i])
        surf.append((gen, dmin, 2 * np.pi * c_r[i] * c_h[i] * np.mean(_weight(gen(k)))))
    g = rng.uniform(-half, half, (200_000, 2))
    w_ground = np.mean(1.0 / np.maximum(np.hypot(g[:, 0], g[:, 1]), R_MIN) ** 2) * (2 * half) ** 2
    wts = np.array([w_ground] + [s[2] for s in surf])
    counts = rng.multinomial(n, wts / wts.sum())

    # Ground: density 1/max(r,1)^2 per area -> uniform disk for r < 1 and
    # log-uniform r beyond, clipped to the square by rejection.
    n_g = int(counts[0])
    r_max = half * np.sqrt(2.0)
    inner_w = np.pi
    outer_w = 2 * np.pi * np.log(r_max / R_MIN)
    parts, have = [], 0
    while have < n_g:
        want = int((n_g - have) * 1.4) + 1024
        inner = rng.random(want) < inner_w / (inner_w + outer_w)
        rr = np.where(inner, R_MIN * np.sqrt(rng.random(want)),
                      R_MIN * np.exp(rng.random(want) * np.log(r_max / R_MIN)))
        th = rng.random(want) * 2 * np.pi
        x, y = rr * np.cos(th), rr * np.sin(th)
        ok = (np.abs(x) <= half) & (np.abs(y) <= half)
        parts.append(np.stack([x[ok], y[ok], np.zeros(int(ok.sum()))], axis=1))
        have += int(ok.sum())
    pieces = [np.concatenate(parts)[:n_g]]
    for (gen, dmin, _), m in zip(surf, counts[1:]):
        if m:
            pieces.append(_sample_weighted(rng, gen, int(m), 1.0 / max(dmin, R_MIN) ** 2))
    pts = np.concatenate(pieces)
    pts += rng.normal(0.0, noise, pts.shape)
    pts = pts[rng.permutation(len(pts))]
    return pts.astype(dtype)


def isotropic_design16():
    """16 points with zero mean and second moment 0.625 I: the 8 cube
    vertices (+-1, +-1, +-1) and the same at half size."""
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    return np.concatenate([v, 0.5 * v])


def random_rotations(rng, m):
    """m rotation matrices, uniform (QR of Gaussian matrices, signs fixed)."""
    a = rng.normal(size=(m, 3, 3))
    q, r = np.linalg.qr(a)
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    det = np.linalg.det(q)
    q[det < 0, :, 0] *= -1
    return q


def _lattice_nearest(count, spacing):
    """The `count` points of a cubic lattice of `spacing` nearest the origin."""
    L = int(np.ceil((3 * count / (4 * np.pi)) ** (1 / 3))) + 2
    g = np.arange(-L, L + 1)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    d = np.sum(pts * pts, axis=1)
    order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0], d))
    return pts[order[:count]] * spacing


def isotropic_groups(n_groups, centre, sigma=0.05, size=4e-4, seed=0):
    """n_groups x 16 points: rotated isotropic designs of half-width `size`
    on the lattice points nearest `centre`, spaced so that the cluster's
    standard deviation per axis is `sigma`; each group jittered by
    N(0, j^2) with j log-uniform in [1e-9, 1e-6] m."""
    rng = np.random.default_rng(seed)
    radius = sigma * np.sqrt(5.0)  # a uniform ball of this radius has std sigma
    spacing = radius / (3 * n_groups / (4 * np.pi)) ** (1 / 3)
    if spacing < 4 * size * np.sqrt(3.0):
        raise ValueError("groups would overlap")
    centres = np.asarray(centre, np.float64) + _lattice_nearest(n_groups, spacing)
    rot = random_rotations(rng, n_groups)
    design = isotropic_design16() * size
    pts = centres[:, None, :] + np.einsum("gij,pj->gpi", rot, design)
    jit = np.exp(rng.uniform(np.log(1e-9), np.log(1e-6), n_groups))
    pts += rng.normal(size=pts.shape) * jit[:, None, None]
    return pts.reshape(-1, 3)


def precision_stress(n_scene=10_000_000, n_cluster=2_000_000, seed=1):
    if n_cluster % 16:
        raise ValueError("n_cluster must be a multiple of 16")
    base = lidar_scene(n_scene, seed=seed, dtype=np.float64)
    cl = isotropic_groups(n_cluster // 16, (10.0, 10.0, 1.0), seed=seed + 100)
    pts = np.concatenate([base, cl]) + np.array([512000.0, 5400000.0, 300.0])
    return pts


def gap_clusters(gaps, per_gap=8, spacing=20.0, seed=0):
    """float32 clusters of 16 points, isolated (`spacing` apart on a grid
    centred at the origin): cluster j of gap g is the isotropic design mapped
    to covariance R diag(1, 1 + 4g, 4) R^T, R a random rotation, so its
    eigengap (l1 - l0) / l2 is g (before float32 rounding, which moves it by
    ~1e-5). Returns (points (n, 3) float32, gap of each point's cluster)."""
    rng = np.random.default_rng(seed)
    gaps = np.repeat(np.asarray(gaps, np.float64), per_gap)
    m = len(gaps)
    side = int(np.ceil(m ** (1 / 3)))
    g = (np.arange(side) - (side - 1) / 2) * spacing
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)[:m]
    rot = random_rotations(rng, m)
    design = isotropic_design16() / np.sqrt(0.625)  # unit second moment
    lam = np.stack([np.ones(m), 1 + 4 * gaps, np.full(m, 4.0)], 1)
    local = design[None, :, :] * np.sqrt(lam)[:, None, :]
    pts = grid[:, None, :] + np.einsum("gij,gpj->gpi", rot, local)
    perm = rng.permutation(m * 16)
    return (pts.reshape(-1, 3)[perm].astype(np.float32), np.repeat(gaps, 16)[perm])


def batched_clouds(count=4096, seed=3, lo=1_000, hi=200_000, max_total=None):
    """List of float32 clouds (sizes log-uniform in [lo, hi])."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
    if max_total is not None:
        keep = np.cumsum(sizes) <= max_total
        sizes = sizes[keep]
    clouds = []
    for s in sizes:
        nprim = rng.integers(1, 6)
        kinds = rng.integers(0, 2, nprim)
        share = rng.multinomial(int(s), np.ones(nprim) / nprim)
        parts = []
        for kind, m in zip(kinds, share):
            if m == 0:
                continue
            if kind == 0:  # plane patch
                c = rng.uniform(0, 10, 3)
                a = rng.normal(size=3)
                a /= np.linalg.norm(a)
                b = np.cross(a, rng.normal(size=3))
                b /= np.linalg.norm(b)
                nrm = np.cross(a, b)
                uv = rng.uniform(-2, 2, (m, 2))
                p = c + uv[:, :1] * a + uv[:, 1:] * b + rng.normal(0, 0.01, (m, 1)) * nrm
            else:  # sphere
                c = rng.uniform(1, 9, 3)
                rad = rng.uniform(0.3, 1.5)
                d = rng.normal(size=(m, 3))
                d /= np.linalg.norm(d, axis=1, keepdims=True)
                p = c + d * (rad + rng.normal(0, 0.01, (m, 1)))
            parts.append(p)
        clouds.append(np.clip(np.concatenate(parts), 0.0, 10.0).astype(np.float32))
    return clouds
