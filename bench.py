"""Benchmark: k-d tree build + kNN + r-near + normals on the 10M-point lidar scene.

One step = one pass of the hot path over the synthetic 10M-point float32
lidar scene (sensor at the origin, ground + 40 walls + 200 cylinders, density
~ 1/range^2, 1 cm noise; pptk_amd.scenes.lidar_scene, seed 1), inputs already
resident in HBM (pptk_amd.pipeline.LidarStep):
  1. build the tree from the device point array;
  2. kNN k=16 self query for every point (idx + counts to HBM);
  3. r-near dmax=0.1 m for 1M query points (sampled without replacement,
     seed 2, queried as points; CSR to HBM);
  4. PCA normals k=16, r=inf for every point on the resident tree.
value = (kNN queries + normals) per second of step time.

Per-phase durations come from HIP events recorded on the stream the library
launches on. The CPU baseline runs pptk's own k-d tree (compiled from the
reference sources, oracle/_ref) or, when that is absent, the oracle's
restatement, on a bounded sample, on the host cores; the same leg checks a
sample of the GPU normals against FP64 normals (angle quantiles). The drop-in
API (host numpy arrays in and out) is timed once, outside the timed steps.

The default run also measures, after the lidar steps, the batched-clouds
workload (4096 clouds, one hip-graph replay per step, bucket 192 and query
block 64: the tuner's choice for it) and the served stream (Poisson
arrivals over a sweep of rates), and reports them as the "batched" and
"served" keys of the same line, each with its ratio to the plain
single-cloud path on the 10M-point cloud (build + kNN + normals).

--workload batched / served runs one of those alone with its full report
(batched: with the autotuner unless --leaf is given).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu]
                       [--no-extras] [--workload lidar|batched|served]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("kNN queries/s + normals/s at 10M pts vs plain GPU & CPU; "
          "served p50/p99 latency")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on (the GPU box's CPU share)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def algorithmic_bytes(n, k, m_rnear, rnear_total):
    s = 4  # float32 coordinates
    return {
        # read points, write reordered coordinates + permutation
        "build": n * (3 * s + 3 * s + 4),
        # read query coordinates, write k int32 indices + 1 int32 count
        "knn": n * (3 * s + 4 * k + 4),
        # read query, write its int32 count and int64 offset
        "rnear_count": m_rnear * (3 * s + 4 + 8),
        # read query and its offsets, write the int32 neighbour ids
        "rnear_fill": m_rnear * (3 * s + 16) + 4 * rnear_total,
        # read point, write the unit normal
        "normals": n * (3 * s + 3 * s),
        # one search: read point, write the kNN row + count and the normal
        "knn_normals": n * (3 * s + 4 * k + 4 + 3 * s),
    }


# The kernel that dominates each phase (names as rocprofv3 reports them).
PHASE_KERNEL = {
    "knn": "knn_coop_kernel<3, 16>",
    "rnear_count": "rnear_packet_kernel<float, 3, false, false>",
    "rnear_fill": "rnear_packet_kernel<float, 3, true, false>",
    "normals": "normals_coop_kernel<16, 1, 1>",
    "knn_normals": "knn_normals_coop_kernel<16, 1, 1>",
}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from profiles/traffic.json (rocprofv3
    FETCH_SIZE and WRITE_SIZE runs scaled by a plain-copy calibration)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)["kernels"][kernel]
        return t["fetch_bytes"] + t["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def quantiles(a):
    a = np.asarray(a)
    return {"p50": float(np.percentile(a, 50)), "p90": float(np.percentile(a, 90)),
            "p99": float(np.percentile(a, 99)), "max": float(np.max(a)), "n": int(a.size)}


# ------------------------------------------------------------ CPU legs ----
# The oracle (oracle/) is imported only here: pptk's CPU path timed as the
# baseline, and FP64 normals of sampled rows as the checker of the GPU's.

def cpu_check_normals(pts, nrm, rows, k, tree=None):
    """Angle (degrees) between GPU normals nrm[rows] and FP64 normals of the
    same neighbourhoods (oracle kNN including the point, LAPACK eigh)."""
    from oracle import oracle as O
    ot = tree if tree is not None else O.OracleTree(pts)
    nb, cnt = ot.knn(pts[rows], k, nthreads=host_cores())
    fev, _ = O.normals_fp64(pts, nb, cnt)
    return O.angle_deg(nrm[rows], fev[:, 0])


def cpu_baseline(pts, qpts, k, dmax, gpu_knn_rows=None, gpu_normals=None):
    """Time pptk's k-d tree (or the oracle port) on a bounded sample, compare
    its kNN rows with the GPU's rows for the same queries, and check a sample
    of the GPU normals against FP64."""
    from oracle import oracle as O
    cores = host_cores()
    n = len(pts)
    kind = "reference" if O.ref_available() else "port"
    t0 = time.perf_counter()
    tree = O.RefTree(pts) if kind == "reference" else O.OracleTree(pts)
    t_build = time.perf_counter() - t0
    mq = min(n, 1_000_000)
    t0 = time.perf_counter()
    if kind == "reference":
        cpu_rows, _ = tree.knn(None, k, nthreads=cores, m_self=mq)
    else:
        cpu_rows, _ = tree.knn(np.arange(mq), k, nthreads=cores)
    t_knn = time.perf_counter() - t0
    same = None
    if gpu_knn_rows is not None:
        same = int(np.sum(np.all(gpu_knn_rows[:mq] == cpu_rows, axis=1)))
    mr = min(len(qpts), 200_000)
    t0 = time.perf_counter()
    tree.rnear(qpts[:mr].astype(pts.dtype), dmax, nthreads=cores)
    t_rn = time.perf_counter() - t0
    del tree
    port = O.OracleTree(pts)
    mn = min(n, 300_000)
    t0 = time.perf_counter()
    port.normals(k, np.inf, rows=np.arange(mn), nthreads=cores)
    t_nrm = time.perf_counter() - t0
    angles = None
    if gpu_normals is not None:
        rows = np.random.default_rng(5).choice(n, size=min(n, 100_000), replace=False)
        angles = quantiles(cpu_check_normals(pts, gpu_normals, rows, k, tree=port))
    del port
    # one host thread, on a smaller sample
    m1 = min(n, 100_000)
    tree1 = O.RefTree(pts) if kind == "reference" else O.OracleTree(pts)
    t0 = time.perf_counter()
    if kind == "reference":
        tree1.knn(None, k, nthreads=1, m_self=m1)
    else:
        tree1.knn(np.arange(m1), k, nthreads=1)
    knn_rate_1t = m1 / (time.perf_counter() - t0)
    del tree1
    knn_rate = mq / t_knn
    nrm_rate = mn / t_nrm
    rn_rate = mr / t_rn
    step = t_build + n / knn_rate + len(qpts) / rn_rate + n / nrm_rate
    base = {
        "value": 2 * n / step,
        "unit": "(kNN queries + normals)/s",
        "cores": cores,
        "kind": kind,
        "sample": ("full %d-point tree build (%s, serial: %.2fs); kNN k=%d self queries on "
                   "%d points; r-near dmax=%g on %d query points; normals k=%d on %d points "
                   "(oracle port, FP64 Jacobi); per-phase rates scaled to the full workload"
                   % (n, "pptk KdTree" if kind == "reference" else "oracle port", t_build, k,
                      mq, dmax, mr, k, mn)),
        "build_s": t_build,
        "knn_qps": knn_rate,
        "normals_ps": nrm_rate,
        "rnear_qps": rn_rate,
        "knn_core_s_full": n / knn_rate,
        "knn_rows_identical_to_cpu": same,
        "knn_rows_compared": mq,
        "knn_qps_1thread": knn_rate_1t,
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
    }
    return base, angles


def cpu_baseline_batched(clouds, k, nrm=None, offsets=None):
    """pptk's CPU path per cloud (kdtree._build + self kNN + normals) on every
    64th cloud: pptk KdTree from the reference sources when built, the oracle
    port for the normals; with the GPU normals, their angle to FP64 on up to
    2,000 rows of each sampled cloud."""
    from oracle import oracle as O
    cores = host_cores()
    kind = "reference" if O.ref_available() else "port"
    sel = list(range(0, len(clouds), 64))
    npts = sum(len(clouds[c]) for c in sel)
    t0 = time.perf_counter()
    for c in sel:
        pts = clouds[c]
        tree = O.RefTree(pts) if kind == "reference" else O.OracleTree(pts)
        if kind == "reference":
            tree.knn(None, k, nthreads=cores, m_self=len(pts))
        else:
            tree.knn(np.arange(len(pts)), k, nthreads=cores)
        del tree
        O.OracleTree(pts).normals(k, np.inf, nthreads=cores)
    t = time.perf_counter() - t0
    base = {
        "value": npts / t,
        "unit": "points/s (kNN k=%d self query + normals k=%d per point)" % (k, k),
        "cores": cores,
        "kind": kind,
        "sample": ("every 64th cloud (%d clouds, %d points): %s build + kNN, oracle-port "
                   "normals (FP64 Jacobi)" % (len(sel), npts,
                                              "pptk KdTree" if kind == "reference"
                                              else "oracle port")),
    }
    angles = None
    if nrm is not None:
        rng = np.random.default_rng(7)
        out = []
        for c in sel:
            pts = clouds[c]
            rows = rng.choice(len(pts), size=min(len(pts), 2000), replace=False)
            out.append(cpu_check_normals(pts, nrm[offsets[c]:offsets[c + 1]], rows, k))
        angles = quantiles(np.concatenate(out))
    return base, angles


def cpu_check_served_normals(pairs, k):
    """Angles of served normals results (cloud points, GPU normals) against
    FP64 on up to 2,000 rows each."""
    rng = np.random.default_rng(8)
    out = []
    for pts, nrm in pairs:
        rows = rng.choice(len(pts), size=min(len(pts), 2000), replace=False)
        out.append(cpu_check_normals(pts, nrm, rows, k))
    return quantiles(np.concatenate(out)) if out else None


# ------------------------------------------------------------ workloads ----

def api_times(pts, k):
    """The drop-in API on host numpy arrays (PCIe transfers and the Python
    result list included), timed once: kdtree._build, kdtree._query(None, k)
    and estimate_normals(k, inf); and the same kNN through the C ABI into one
    host array, whose difference to _query is the list construction."""
    import pptk_amd
    from pptk_amd import capi
    t0 = time.perf_counter()
    tree = pptk_amd.kdtree._build(pts)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = pptk_amd.kdtree._query(tree, None, k=k)
    t_query = time.perf_counter() - t0
    assert len(res) == len(pts)
    t0 = time.perf_counter()
    del res
    t_free = time.perf_counter() - t0
    del tree
    ct = capi.Tree.build(pts)
    t0 = time.perf_counter()
    idx, cnt = capi.knn(ct, None, k)
    t_core = time.perf_counter() - t0
    ct.free()
    del idx, cnt
    t0 = time.perf_counter()
    nrm = pptk_amd.estimate_normals(pts, k, np.inf, verbose=False)
    t_nrm = time.perf_counter() - t0
    del nrm
    return {
        "build_s": t_build,
        "query_knn_s": t_query,
        "query_knn_one_array_s": t_core,
        "query_knn_list_s": max(0.0, t_query - t_core),
        "result_list_free_s": t_free,
        "estimate_normals_s": t_nrm,
        "what": ("drop-in calls on host numpy arrays: kdtree._build(pts); "
                 "kdtree._query(tree, None, k=%d) -> list of %d int32 arrays; "
                 "estimate_normals(pts, %d, inf) (its own tree); query_knn_list_s = "
                 "_query minus the same kNN into one host array" % (k, len(pts), k)),
    }


def bench_lidar(args):
    import torch

    from pptk_amd import capi, pipeline, scenes

    n, k, dmax = args.n, args.k, float(np.float32(args.dmax))
    log("bench: generating lidar scene n=%d" % n)
    pts = scenes.lidar_scene(n, seed=1)
    rng = np.random.default_rng(2)
    qsel = rng.choice(n, size=min(args.queries, n), replace=False)
    qpts = np.ascontiguousarray(pts[qsel])
    m = len(qpts)

    dev = torch.device("cuda:0")
    pts_d = torch.from_numpy(pts).to(dev)
    q_d = torch.from_numpy(qpts).to(dev)
    prec = capi.PREC_MIXED if args.precision == "mixed" else capi.PREC_FP64
    pipe = pipeline.LidarStep(pts_d, q_d, k=k, dmax=dmax, leaf=args.leaf, precision=prec)

    def timed(p, warmup, steps, say):
        """(ms per step by the host clock, mean phase ms by events)."""
        phases = list(p.PHASES)
        rec = {}
        for _ in range(warmup):
            p.run(None)
        # the phase events are created before and read after the timed steps
        # (every library call returns with its stream synchronised)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(len(phases) + 1)]
               for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            p.run(evs[i])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        for i, ev in enumerate(evs):
            for j, name in enumerate(phases):
                rec.setdefault(name, []).append(ev[j].elapsed_time(ev[j + 1]))
            if say:
                log("bench: step %d: %.3f ms of phases" % (i, ev[0].elapsed_time(ev[-1])))
        return wall * 1e3 / steps, {q: float(np.mean(v)) for q, v in rec.items()}

    log("bench: warmup %d, timing %d steps" % (args.warmup, args.steps))
    ms_step, ph = timed(pipe, args.warmup, args.steps, True)
    sep_ms, sep_ph = None, None
    if not args.only_step:
        # the same step with the two separate calls (two searches per point)
        log("bench: the step with separate kNN and normals calls")
        sep = pipeline.LidarStep(pts_d, q_d, k=k, dmax=dmax, leaf=args.leaf, precision=prec,
                                 fused=False)
        sep_ms, sep_ph = timed(sep, 2, min(args.steps, 5), False)
        del sep

    ab = algorithmic_bytes(n, k, m, pipe.rtotal)
    dom = max((p for p in PHASE_KERNEL if p in ph), key=lambda p: ph[p])
    achieved = ab[dom] / (ph[dom] * 1e-3) / 1e9
    traffic = measured_traffic(PHASE_KERNEL[dom])
    # the single-cloud path for kNN + normals of every point of the cloud
    plain_ms = ph["build"] + ph["knn_normals"]
    out = {
        "metric": METRIC,
        "value": 2 * n / (ms_step * 1e-3),
        "unit": "(kNN queries + normals)/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("lidar scene %d float32 points: build + kNN k=%d self query + r-near "
                         "dmax=%g m on %d query points + normals k=%d r=inf (%s)"
                         % (n, k, args.dmax, m, k, args.precision)),
            "n_points": n, "k": k, "rnear_queries": m, "dmax": args.dmax,
            "leaf_size": pipe.leaf_size,
        },
        "phases_ms": ph,
        "knn_normals_points_per_s": n / (ph["knn_normals"] * 1e-3),
        "separate_calls": None if sep_ph is None else {
            "what": "the same step with pptk_knn then pptk_normals (a search each) in place "
                    "of pptk_knn_normals; outputs bit-identical",
            "ms_per_step": sep_ms, "phases_ms": sep_ph,
            "knn_qps": n / (sep_ph["knn"] * 1e-3),
            "normals_ps": n / (sep_ph["normals"] * 1e-3),
        },
        "rnear_qps": m / ((ph["rnear_count"] + ph["rnear_fill"]) * 1e-3),
        "build_pts_per_s": n / (ph["build"] * 1e-3),
        "plain_10m_points_per_s": n / (plain_ms * 1e-3),
        "rnear_neighbours": pipe.rtotal,
        "normals_fp64_fallback_frac": pipe.num_fp64 / n,
        "roofline": {
            "bound": "hbm",
            "kernel": PHASE_KERNEL[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": ab[dom],
        },
    }
    rows = pipe.idx_d[: min(n, 1_000_000)].cpu().numpy()
    nrm = pipe.nrm_d.cpu().numpy()
    del pipe, pts_d, q_d
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    capi.lib().pptk_trim()
    if args.only_step:
        return out
    log("bench: drop-in API timing")
    out["api"] = api_times(pts, k)
    if not args.no_cpu:
        log("bench: CPU baseline")
        base, angles = cpu_baseline(pts, qpts, k, dmax, gpu_knn_rows=rows, gpu_normals=nrm)
        out["cpu_baseline"] = base
        out["normals_angle_deg_vs_fp64"] = angles
        out["api"]["cpu_knn_api_estimate_s"] = (base["knn_core_s_full"] +
                                                out["api"]["query_knn_list_s"])
    return out


BATCH_KERNEL = "batch_query_kernel<float, 3, 16, 1, true>"


def bench_batched(args, plain_10m_pps=None):
    """Batched clouds: one plan over all clouds (forest build + fused kNN and
    normals, one hip-graph replay per step) against the plain single-cloud
    path (pptk_tree_build + pptk_knn + pptk_normals per cloud) on the same
    clouds."""
    import torch

    from pptk_amd import capi, scenes, tune

    k = args.k
    log("bench: generating %d batched clouds (seed 3)" % args.clouds)
    clouds = scenes.batched_clouds(args.clouds, seed=3)
    sizes = np.array([len(c) for c in clouds], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)])
    N = int(off[-1])
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    sh = stream.cuda_stream
    pts_d = torch.from_numpy(np.concatenate(clouds)).to(dev)
    idx_d = torch.empty((N, k), dtype=torch.int32, device=dev)
    nrm_d = torch.empty((N, 3), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    if args.leaf:
        best, table = {"leaf": args.leaf, "block": args.qblock, "ms": None}, []
    else:
        # autotune on every 8th cloud
        sel = np.arange(0, len(sizes), 8)
        sample_d = torch.cat([pts_d[off[c]:off[c + 1]] for c in sel])
        torch.cuda.synchronize()
        log("bench: autotuning on %d clouds (%d points)" % (len(sel), sample_d.shape[0]))
        best, table = tune.tune_batch(sample_d, sizes[sel], k=k, normals_k=k, stream=stream,
                                      log=log, use_cache=False)
        del sample_d
    log("bench: leaf %d block %d" % (best["leaf"], best["block"]))

    plan = capi.Batch(sizes, k=k, normals_k=k, leaf=best["leaf"], query_block=best["block"],
                      use_graph=True)
    run_args = (pts_d.data_ptr(), idx_d.data_ptr(), None, nrm_d.data_ptr(), None)
    for _ in range(args.warmup):
        plan.run(*run_args, stream=sh)
    # events created before and read after the timed replays
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        evs[i][0].record(stream)
        plan.run(*run_args, stream=sh)
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    times = [e0.elapsed_time(e1) for e0, e1 in evs]
    for i, t in enumerate(times):
        log("bench: batched step %d %.2f ms" % (i, t))
    binfo = plan.info()

    # phase split from one eager timed run (same kernels, no graph)
    tplan = capi.Batch(sizes, k=k, normals_k=k, leaf=best["leaf"], query_block=best["block"],
                       use_graph=False, timing=True)
    tplan.run(*run_args, stream=sh)
    tinfo = tplan.info()
    tplan.free()

    if args.no_plain:
        plan.free()
        return {"workload": "batched", "ms_per_step": ms_step,
                "phases_ms_eager": {"build": tinfo.build_ms, "query": tinfo.query_ms}}
    # plain single-cloud path on the same clouds
    log("bench: plain single-cloud path over %d clouds" % len(sizes))
    L = capi.lib()
    pidx_d = torch.empty((N, k), dtype=torch.int32, device=dev)
    pcnt_d = torch.empty(N, dtype=torch.int32, device=dev)
    pnrm_d = torch.empty((N, 3), dtype=torch.float32, device=dev)
    qall, _ = capi.make_queries(None, mem=capi.MEM_DEVICE)
    qall.kind = capi.Q_ALL
    shp = ctypes.c_void_p(sh)

    def plain(c):
        a, n = int(off[c]), int(sizes[c])
        tree = capi.Tree.build(dev_ptr=pts_d[a:].data_ptr(), n=n, dim=3, dtype=np.float32,
                               leaf=0, stream=shp)
        capi.check(L.pptk_knn(ctypes.c_void_p(tree.handle), ctypes.byref(qall), k,
                              float("inf"), ctypes.c_void_p(pidx_d[a:].data_ptr()),
                              ctypes.c_void_p(pcnt_d[a:].data_ptr()), None, capi.MEM_DEVICE,
                              shp))
        capi.check(L.pptk_normals(ctypes.c_void_p(tree.handle), k, float("inf"), None, 0, 1,
                                  capi.PREC_MIXED, ctypes.c_void_p(pnrm_d[a:].data_ptr()), None,
                                  None, None, capi.MEM_DEVICE, shp))
        tree.free()

    for c in range(min(8, len(sizes))):  # warm the allocator
        plain(c)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in range(len(sizes)):
        plain(c)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    same_idx = bool(torch.equal(pidx_d, idx_d))
    same_nrm = bool(torch.equal(pnrm_d, nrm_d))
    del pidx_d, pcnt_d, pnrm_d

    value = N / (ms_step * 1e-3)
    plain_pps = N / t_plain
    ab = N * (3 * 4 + 4 * k + 3 * 4)
    achieved = ab / (tinfo.query_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s (kNN k=%d self query + normals k=%d per point, forest build "
                "included)" % (k, k),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("%d batched float32 clouds, sizes log-uniform [1e3, 2e5] (seed 3), "
                         "%d points: forest build + kNN k=%d self query + normals k=%d, one "
                         "hip-graph replay per step" % (len(sizes), N, k, k)),
            "clouds": int(len(sizes)), "n_points": N, "k": k,
            "leaf_size": best["leaf"], "query_block": best["block"],
        },
        "graph_ms_per_step": float(np.mean(times)),
        "phases_ms_eager": {"build": tinfo.build_ms, "query": tinfo.query_ms},
        "plain_path": {
            "points_per_s": plain_pps,
            "seconds": t_plain,
            "what": "per cloud: pptk_tree_build + pptk_knn(all) + pptk_normals, same stream",
            "results_identical_to_batched": {"knn": same_idx, "normals": same_nrm},
        },
        "speedup_vs_plain_path_per_cloud": value / plain_pps,
        "speedup_vs_plain_path_10m_cloud": (value / plain_10m_pps if plain_10m_pps
                                            else None),
        "batch": {"num_nodes": binfo.num_nodes, "depth": binfo.depth,
                  "morton_bits": binfo.morton_bits, "kernel_launches": binfo.graph_nodes,
                  "device_bytes": binfo.device_bytes,
                  "normals_fp64_fallback_frac": binfo.num_fp64 / N},
        "tuning": table,
        "roofline": {
            "bound": "hbm",
            "kernel": BATCH_KERNEL,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(BATCH_KERNEL),
            "algorithmic_bytes_per_launch": ab,
        },
    }
    plan.free()
    nrm = nrm_d.cpu().numpy()
    del pts_d, idx_d, nrm_d
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    capi.lib().pptk_trim()
    if not args.no_cpu:
        log("bench: CPU baseline (batched)")
        base, angles = cpu_baseline_batched(clouds, k, nrm=nrm, offsets=off)
        out["cpu_baseline"] = base
        out["normals_angle_deg_vs_fp64"] = angles
    return out


def bench_served(args, plain_10m_pps=None):
    """Served stream: Poisson arrivals of mixed requests over lidar clouds of
    log-uniform size [1e4, 2e6] (pptk_amd.serve.request_mix), swept over
    offered rates, against the plain path (the same requests one at a time,
    a tree built per request)."""
    import torch

    from pptk_amd import capi, scenes, serve

    torch.cuda.is_available()
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(2e6), args.pool)).astype(np.int64)
    log("bench: generating %d lidar clouds (%d points)" % (args.pool, sizes.sum()))
    clouds = [scenes.lidar_scene(int(n), seed=1000 + i) for i, n in enumerate(sizes)]
    rates = [float(x) for x in args.rates.split(",")]
    R = args.requests

    def sequence():
        return serve.request_mix(np.random.default_rng(5), clouds, R)

    # warm the allocator, the library and torch's streams
    w = serve.Server(serve.GpuBackend(args.slots, outputs=args.outputs), slots=args.slots)
    for t in [w.submit(r) for r in sequence()[:16]]:
        t.result()
    w.close()

    sweep = []
    normals_pairs = []
    for rate in rates:
        server = serve.Server(serve.GpuBackend(args.slots, outputs=args.outputs),
                              slots=args.slots, budget_bytes=200 * 10 ** 9)
        tickets, t0, t1 = serve.run_open_loop(server, sequence(), rate,
                                              np.random.default_rng(6))
        st = serve.percentile_stats(tickets, t0, t1)
        st["offered_rate"] = rate
        st["peak_reserved_plus_resident_bytes"] = server.memory()["peak"]
        st["trees_built"] = sum(1 for t in tickets if t.built)
        if rate == rates[-1]:
            for t in tickets:
                if t.req.kind == "normals" and t.error is None and len(normals_pairs) < 4:
                    res = t.result()
                    res = res.cpu().numpy() if hasattr(res, "cpu") else np.asarray(res)
                    normals_pairs.append((clouds[int(t.req.cloud[5:]) % len(clouds)], res))
        server.close()
        sweep.append(st)
        log("bench: rate %g req/s -> %.1f req/s, %.3g pts/s, p50 %.1f ms, p99 %.1f ms"
            % (rate, st["requests_per_s"], st["points_per_s"], st["p50_ms"], st["p99_ms"]))

    # plain path: each request alone, its cloud's tree built for it
    log("bench: plain path, %d requests one at a time" % R)
    be = serve.GpuBackend(1, outputs=args.outputs)
    lat = []
    npts = 0
    t0 = time.perf_counter()
    for r in sequence():
        ts = time.perf_counter()
        pts = clouds[int(r.cloud[5:]) % len(clouds)]
        tree = be.build(pts, 0)
        res = be.run(r, tree, 0, 1 << 30)
        del res
        tree.free()
        npts += r.size
        lat.append(time.perf_counter() - ts)
    t_plain = time.perf_counter() - t0
    plain = {"points_per_s": npts / t_plain, "requests_per_s": R / t_plain,
             "p50_ms": float(np.percentile(lat, 50) * 1e3),
             "p99_ms": float(np.percentile(lat, 99) * 1e3),
             "what": "same request sequence, one at a time, build + query per request"}
    base_p99 = sweep[0]["p99_ms"]
    ok = [st for st in sweep if st["requests_per_s"] >= 0.9 * st["offered_rate"]
          and st["p99_ms"] <= 3.0 * base_p99]
    knee = ok[-1] if ok else sweep[0]
    best = max(sweep, key=lambda st: st["points_per_s"])
    out = {
        "metric": METRIC,
        "value": knee["points_per_s"],
        "unit": "queried points/s served at the knee rate",
        "n_gpus": 1,
        "steps": R,
        "warmup": 16,
        "ms_per_step": 1e3 / knee["requests_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("served stream: %d requests per rate, Poisson arrivals, lidar clouds "
                         "log-uniform [1e4, 2e6] points (%d-cloud pool), kinds uniform over "
                         "kNN k=8/16/32, r-near dmax=0.05/0.2 m (20,000 self queries), "
                         "normals k=16; half follow-ups on resident trees; %d slots; outputs "
                         "in %s memory; a request counts the points it queries"
                         % (R, args.pool, args.slots, args.outputs)),
            "slots": args.slots, "hbm_budget_bytes": 200 * 10 ** 9, "outputs": args.outputs,
        },
        "knee": {"offered_rate": knee["offered_rate"], "p50_ms": knee["p50_ms"],
                 "p99_ms": knee["p99_ms"], "requests_per_s": knee["requests_per_s"],
                 "points_per_s": knee["points_per_s"]},
        "sweep": sweep,
        "plain_path": plain,
        "speedup_vs_plain_path_at_knee": knee["points_per_s"] / plain["points_per_s"],
        "speedup_vs_plain_path_best": best["points_per_s"] / plain["points_per_s"],
        "speedup_vs_plain_path_10m_cloud_at_knee": (knee["points_per_s"] / plain_10m_pps
                                                    if plain_10m_pps else None),
    }
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    capi.lib().pptk_trim()
    if not args.no_cpu:
        out["normals_angle_deg_vs_fp64"] = cpu_check_served_normals(normals_pairs, 16)
    return out


def summary_batched(b):
    keep = ("value", "unit", "ms_per_step", "config", "phases_ms_eager",
            "speedup_vs_plain_path_per_cloud", "speedup_vs_plain_path_10m_cloud",
            "roofline", "cpu_baseline", "normals_angle_deg_vs_fp64")
    out = {k: b[k] for k in keep if k in b}
    out["plain_path"] = {k: b["plain_path"][k] for k in ("points_per_s",
                                                         "results_identical_to_batched")}
    out["normals_fp64_fallback_frac"] = b["batch"]["normals_fp64_fallback_frac"]
    return out


def summary_served(s):
    keep = ("value", "unit", "config", "knee", "plain_path", "speedup_vs_plain_path_at_knee",
            "speedup_vs_plain_path_best", "speedup_vs_plain_path_10m_cloud_at_knee",
            "normals_angle_deg_vs_fp64")
    out = {k: s[k] for k in keep if k in s}
    out["sweep"] = [{k: st[k] for k in ("offered_rate", "requests_per_s", "points_per_s",
                                        "p50_ms", "p99_ms")} for st in s["sweep"]]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="lidar", choices=["lidar", "batched", "served"])
    ap.add_argument("--rates", default="400,800,1200,1600,2000,2400,2800,3200,4000",
                    help="served: offered request rates (requests/s)")
    ap.add_argument("--requests", type=int, default=200, help="served: requests per rate")
    ap.add_argument("--pool", type=int, default=40, help="served: distinct clouds")
    ap.add_argument("--slots", type=int, default=16, help="served: concurrent requests")
    ap.add_argument("--outputs", default="device", choices=["device", "host"],
                    help="served: where results are returned")
    ap.add_argument("--clouds", type=int, default=4096)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--dmax", type=float, default=0.1)
    ap.add_argument("--leaf", type=int, default=0)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--qblock", type=int, default=64,
                    help="batched: query block width when --leaf fixes the bucket size")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="lidar: skip the batched and served summaries")
    ap.add_argument("--only-step", action="store_true",
                    help="lidar: only the timed step (no separate-call step, API timing, CPU "
                         "baseline or extras; profiling runs)")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched: skip the plain-path comparison (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (counter calibration)")
    ap.add_argument("--calibrate-scan", type=int, default=0, metavar="MIB",
                    help="only run the library's exclusive scan over MIB mebibytes of int32 "
                         "(reads its input twice, writes int64 offsets; counter calibration)")
    ap.add_argument("--calibrate-gather", type=int, default=0, metavar="M",
                    help="only gather M float32 elements from M distinct 128-byte lines of a "
                         "1 GiB array (counter calibration of line granularity)")
    args = ap.parse_args()
    if args.calibrate_scan:
        import torch

        from pptk_amd import capi
        n = (args.calibrate_scan << 20) // 4
        x = torch.ones(n, dtype=torch.int32, device="cuda:0")
        out = torch.empty(n + 1, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        for _ in range(3):
            capi.check(capi.lib().pptk_exclusive_scan_i32(
                ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), n,
                capi.MEM_DEVICE))
        torch.cuda.synchronize()
        assert int(out[-1]) == n
        print(json.dumps({"scan_elements": n, "bytes_read_per_scan": 8 * n,
                          "bytes_written_per_scan": 8 * (n + 1), "scans": 3}), flush=True)
        return
    if args.calibrate_gather:
        import torch
        m = args.calibrate_gather
        lines = (1 << 30) // 128
        if m > lines:
            raise SystemExit("at most %d lines" % lines)
        src = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda:0")
        g = torch.Generator(device="cpu").manual_seed(9)
        line = torch.randperm(lines, generator=g)[:m]
        idx = (line * 32 + torch.randint(0, 32, (m,), generator=g)).to("cuda:0")
        torch.cuda.synchronize()
        for _ in range(3):
            out = torch.index_select(src, 0, idx)
        torch.cuda.synchronize()
        print(json.dumps({"gathered": m, "distinct_lines_128B": m, "index_bytes": 8 * m,
                          "out_bytes": 4 * m, "gathers": 3, "sum": float(out.sum())}),
              flush=True)
        return
    if args.calibrate_copy:
        import torch
        nbytes = args.calibrate_copy << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported (queries would shard "
              "by replicating the tree, batched clouds and served clouds by cloud)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "served":
        print(json.dumps(bench_served(args)), flush=True)
        return
    out = bench_lidar(args)
    if not (args.no_extras or args.only_step):
        plain10 = out["plain_10m_points_per_s"]
        ex = argparse.Namespace(**vars(args))
        ex.k, ex.leaf, ex.qblock, ex.no_plain = 16, 192, 64, False
        out["batched"] = summary_batched(bench_batched(ex, plain10))
        out["served"] = summary_served(bench_served(ex, plain10))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
